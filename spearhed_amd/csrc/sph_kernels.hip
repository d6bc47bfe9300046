// sph_kernels.hip -- hand-written gfx950 kernels of the SPH-MHD particle step.
//
// Stage -> kernels (reference stage in brackets):
//   drift + bin   k_drift_bin        [PushVelocity, PushVelocity.hpp:35-42; + supercell key/histogram]
//   resort        k_scan_local/top/add, k_scatter, k_gather
//                                    [UpdateVolumes RegionBoundsUpdate.hpp:175-214 and the frame index
//                                     FrameIndex.hpp:209-296, replaced by a counting sort by (cell, id)]
//   density       k_density<CH>      [UpdateDensity, DensitySummation.hpp:104-138]
//   forces        k_forces<CH, MHD>  [UpdateHydroForces, HydroForces.hpp:183-214, + AV/MHD/cleaning]
//   integrate     k_kick, k_finalize [EulerIntegrate, EulerIntegrate.hpp:43-52, + CFL and FP64 sums]
//
// Pair passes: one wave64 per target supercell. The 27 neighbour supercells form a virtual list
// whose candidates within the support radius of the targets' bounding box are staged chunk by
// chunk into LDS (records pre-shifted by the periodic image or the supercell-origin offset,
// ParticleParticleInteraction.hpp:236-264). The cell's nt targets are
// spread over the wave as (target t, sub-lane k) with K = floor(64/nt) sub-lanes per target, so a
// supercell of ~14 particles still fills the wave; sub-lane k owns candidates v = c*K + k.
// The density pass culls each lane's candidates 32 at a time into a bit mask (strict r2 < H^2,
// ParticleParticleInteraction.hpp:291) and stores the masks; the force pass replays them instead
// of culling again. Each lane then runs the pair physics over its set bits; with pair_balance the
// accepted candidates of a target's K lanes are first compacted into LDS and shared round-robin
// among those K lanes (ceil(accepted/K) iterations instead of the slowest lane's count, at the
// price of a prefix scan and two barriers per 32 candidates). The K partial sums are combined in lane order:
// every summation order depends only on the run's own data, so a run gives bit-identical results
// alone or batched with others.

#include <type_traits>

#include "sph_device.hpp"
#include "sph_launch.hpp"

namespace sph {

// ------------------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int find_run_by_cell(const RunDev* runs, int nruns, int cell)
{
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (runs[mid].c0 <= cell) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// An LDS record at a byte offset from its array: the pair loops keep byte offsets so that a slot's
// address is one 24-bit multiply-add (candidate bit x lane stride + the lane's first slot).
template <class T>
__device__ __forceinline__ const T& lds_at(const T* base, int byte_off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ int wave_incl_max(int v)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v = v > t ? v : t;
    }
    return v;
}

// XCD-aware mapping: the hardware deals consecutive workgroups round-robin to the 8 XCDs; give
// each XCD one contiguous range of supercells so that neighbouring cells share an L2.
__device__ __forceinline__ int xcd_cell(int bid, int nblk_total, int ncell)
{
    int per = nblk_total >> 3;
    int cell = (bid & 7) * per + (bid >> 3);
    return cell < ncell ? cell : -1;
}

// ------------------------------------------------------------------------------------------------
// drift + supercell key + histogram rank (wave-aggregated atomics: particles arrive sorted by
// cell, so runs of equal keys in a wave take one atomicAdd per run)
// ------------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(256) void k_drift_bin(const RunDev* __restrict__ runs, StateBuf S, WorkBuf W,
                                                   double dt_host, int do_drift)
{
    const int b = blockIdx.x;
    const int r = W.blk_run[b];
    const RunDev& R = runs[r];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    const bool active = local < R.np;
    const int64_t i = R.p0 + (active ? local : 0);
    uint32_t key = 0xffffffffu;
    if (active) {
        const double dt = dt_host > 0.0 ? dt_host : W.dt[r];
        const uint8_t live = S.mask[i];
        // a run whose next step primes it (fresh initial data) does not move in that step, nor does
        // a run without a CFL bound (no live target: dt = +inf)
        const bool mov = do_drift && is_movable(S.species[i], live) && !W.prime_flag[r] && isfinite(dt);
        int c[3];
        for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
            int ci;
            if (MIXED) {
                double x = S.pos64[a][i];
                if (mov) {
                    x = x + (double)S.vel[a][i] * dt;
                    if (R.periodic[a]) {
                        double hi = R.lo[a] + R.L[a];
                        if (x >= hi) x = x - R.L[a];
                        else if (x < R.lo[a]) x = x + R.L[a];
                    }
                    S.pos64[a][i] = x;
                }
                ci = (int)fmin(fmax(floor((x - R.lo[a]) * R.inv_s[a]), 0.0), (double)(R.n[a] - 1));
            } else {
                float x = S.pos32[a][i];
                if (mov) {
                    x = x + S.vel[a][i] * (float)dt;
                    if (R.periodic[a]) {
                        float hi = R.lo32[a] + R.L32[a];
                        if (x >= hi) x = x - R.L32[a];
                        else if (x < R.lo32[a]) x = x + R.L32[a];
                    }
                    S.pos32[a][i] = x;
                }
                ci = (int)fminf(fmaxf(floorf((x - R.lo32[a]) * R.inv_s32[a]), 0.0f), (float)(R.n[a] - 1));
            }
            // clamped before the conversion: outside the box (or not finite) maps to an edge cell
            c[a] = ci;
        }
        key = (uint32_t)(R.c0 + c[0] + R.n[0] * (c[1] + R.n[1] * c[2]));
        W.key[i] = key;
    }
    // segments of equal consecutive keys within the wave
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prev != key || !active;
    const int hl = wave_incl_max(head ? lane : 0);
    const uint64_t heads = __ballot(head);
    uint32_t base = 0;
    if (head && active) {
        const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
        const int next = after ? __builtin_ctzll(after) : 64;
        base = atomicAdd(&W.count[key], (uint32_t)(next - lane));
    }
    base = __shfl(base, hl, 64);
    if (active) W.rank[i] = base + (uint32_t)(lane - hl);
}

// ------------------------------------------------------------------------------------------------
// exclusive scan of the per-cell histogram -> cell starts (absolute particle index)
// ------------------------------------------------------------------------------------------------
constexpr int kScanPerThread = 8;
constexpr int kScanBlock = 256 * kScanPerThread;

__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* __restrict__ count, uint32_t* __restrict__ cs,
                                                    uint32_t* __restrict__ bsum, int n)
{
    __shared__ uint32_t wsum[4];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int e = base + k;
        v[k] = e < n ? count[e] : 0u;
        s += v[k];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = wave_incl_scan((int)s);
    if (lane == 63) wsum[wid] = (uint32_t)incl;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wid; ++w) woff += wsum[w];
    uint32_t run = woff + (uint32_t)incl - s;
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k) {
        int e = base + k;
        if (e < n) cs[e] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = woff + (uint32_t)incl;
}

__global__ __launch_bounds__(1024) void k_scan_top(uint32_t* __restrict__ bsum, int nb, uint32_t* __restrict__ cs,
                                                   int n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int e = base + threadIdx.x;
        uint32_t s = e < nb ? bsum[e] : 0u;
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        int incl = wave_incl_scan((int)s);
        if (lane == 63) wsum[wid] = (uint32_t)incl;
        __syncthreads();
        uint32_t woff = carry;
        for (int w = 0; w < wid; ++w) woff += wsum[w];
        if (e < nb) bsum[e] = woff + (uint32_t)incl - s;
        __syncthreads();
        if (threadIdx.x == 1023) carry = woff + (uint32_t)incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) cs[n] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ cs, const uint32_t* __restrict__ bsum, int n)
{
    const int base = blockIdx.x * kScanBlock;
    const uint32_t off = bsum[blockIdx.x];
    for (int k = threadIdx.x; k < kScanBlock; k += 256) {
        int e = base + k;
        if (e < n) cs[e] += off;
    }
}

// ------------------------------------------------------------------------------------------------
// scatter (histogram slot) and gather (final slot = cell start + rank of id within the cell)
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(const RunDev* __restrict__ runs, StateBuf S, WorkBuf W)
{
    const int b = blockIdx.x;
    const RunDev& R = runs[W.blk_run[b]];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    if (local >= R.np) return;
    const int64_t i = R.p0 + local;
    const uint32_t slot = W.cs[W.key[i]] + W.rank[i];
    W.tmp[slot] = (uint32_t)i;
    W.tmpid[slot] = S.id[i];
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_gather(const RunDev* __restrict__ runs, StateBuf Sold, StateBuf Snew,
                                                WorkBuf W, int has_mhd)
{
    const int b = blockIdx.x;
    const int r = W.blk_run[b];
    const RunDev& R = runs[r];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    if (local >= R.np) return;
    const int64_t q = R.p0 + local;
    if (local == 0) W.chbits[r] = 0u;  // cleaning-speed maximum of the coming density pass
    const uint32_t i = W.tmp[q];
    const uint32_t c = W.key[i];
    const uint64_t myid = W.tmpid[q];
    // the particle's old record is read before the rank search, so its loads overlap the search's
    const uint8_t live = Sold.mask[i], sp = Sold.species[i];
    const float m = Sold.mass[i], hh = Sold.h[i], uu = Sold.u[i], rh = Sold.rho[i];
    float vel[3], bf[3] = {0.f, 0.f, 0.f}, ps = 0.f;
    double x64[3] = {0.0, 0.0, 0.0};
    float x32[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < 3; ++a) {
        vel[a] = Sold.vel[a][i];
        if (MIXED) x64[a] = Sold.pos64[a][i];
        else x32[a] = Sold.pos32[a][i];
    }
    if (has_mhd) {
        for (int a = 0; a < 3; ++a) bf[a] = Sold.B[a][i];
        ps = Sold.psi[i];
    }
    const uint32_t cb = W.cs[c], ce = W.cs[c + 1];
    // rank within the supercell = number of smaller ids in it; four independent loads per round
    uint32_t rank = 0, mm = cb;
    for (; mm + 4 <= ce; mm += 4) {
        const uint64_t a0 = W.tmpid[mm], a1 = W.tmpid[mm + 1], a2 = W.tmpid[mm + 2], a3 = W.tmpid[mm + 3];
        rank += (a0 < myid ? 1u : 0u) + (a1 < myid ? 1u : 0u) + (a2 < myid ? 1u : 0u) + (a3 < myid ? 1u : 0u);
    }
    for (; mm < ce; ++mm) rank += (W.tmpid[mm] < myid) ? 1u : 0u;
    const uint32_t f = cb + rank;
    if (rank == 0) W.count[c] = 0u;  // histogram reset for the next resort
    if (W.perm) W.perm[f] = (int32_t)i;

    Snew.id[f] = myid;
    Snew.mask[f] = live;
    Snew.species[f] = sp;
    Snew.mass[f] = m;
    Snew.h[f] = hh;
    Snew.u[f] = uu;
    Snew.rho[f] = rh;
    float pc[3];
    const int lc = (int)c - R.c0;
    const int cxyz[3] = {lc % R.n[0], (lc / R.n[0]) % R.n[1], lc / (R.n[0] * R.n[1])};
    for (int a = 0; a < 3; ++a) {
#pragma clang fp contract(off)
        Snew.vel[a][f] = vel[a];
        if (MIXED) {
            const double x = x64[a];
            Snew.pos64[a][f] = x;
            pc[a] = (float)((x - R.lo[a]) - (double)cxyz[a] * R.s[a]);
        } else {
            Snew.pos32[a][f] = x32[a];
            pc[a] = x32[a];
        }
    }
    if (has_mhd) {
        for (int a = 0; a < 3; ++a) Snew.B[a][f] = bf[a];
        Snew.psi[f] = ps;
    }
    float4 p0 = is_source(sp, live) ? make_float4(pc[0], pc[1], pc[2], m)
                                    : make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
    W.P0[f] = p0;
}

// ------------------------------------------------------------------------------------------------
// shared set-up of the supercell stencil's virtual neighbour list (one wave): the (2 reach + 1)^3
// supercells around a cell, x fastest, concatenated. NS = 27 serves supercells of edge >= H (reach
// 1), NS = 125 also those of edge >= H/2 (reach 2: 5^3 stencil, about 58% of the candidates).
// ------------------------------------------------------------------------------------------------
template <int NS>
struct NbListT {
    int start[NS];
    int off[NS + 1];  // entries past the run's stencil hold the total, so no search selects them
    float sh[NS][3];
};
template <int RMAX>
using NbListR = NbListT<RMAX == 1 ? 27 : 125>;
// LDS candidate slot indices of the balanced pair loop
template <int CH>
using ListT = typename std::conditional<(CH <= 256), uint8_t, uint16_t>::type;

// Neighbour supercell `seg` of cell c in a stencil of side 2 reach + 1 (x fastest, offsets
// d = seg % side - reach, ...): returns false outside a non-periodic axis; else the global cell and
// the staged-position shift (the periodic image for FP32 absolute coordinates, d * cell edge for
// supercell-relative ones). A 1-cell periodic axis is wrapped twice by a reach of 2.
template <bool MIXED, int RMAX>
__device__ __forceinline__ bool nb_cell(const RunDev& R, const int c[3], int seg, int& cc, float sh[3])
{
    int d[3];
    if (RMAX == 1) {
        d[0] = seg % 3 - 1;
        d[1] = (seg / 3) % 3 - 1;
        d[2] = seg / 9 - 1;
    } else {
        const int r = R.reach, side = 2 * R.reach + 1;
        d[0] = seg % side - r;
        d[1] = (seg / side) % side - r;
        d[2] = seg / (side * side) - r;
    }
    int idx[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        int x = c[a] + d[a];
        int wrap = 0;
        if (R.periodic[a]) {
            if (RMAX == 1) wrap = x < 0 ? -1 : (x >= R.n[a] ? 1 : 0);
            else wrap = x >= 0 ? x / R.n[a] : -((R.n[a] - 1 - x) / R.n[a]);
            x -= wrap * R.n[a];
        } else if (x < 0 || x >= R.n[a]) {
            ok = false;
        }
        idx[a] = x;
        sh[a] = MIXED ? (float)d[a] * R.s32[a] : (float)wrap * R.L32[a];
    }
    cc = R.c0 + idx[0] + R.n[0] * (idx[1] + R.n[1] * idx[2]);
    return ok;
}

// Index of the cell's own supercell in its stencil (the centre).
template <int RMAX>
__device__ __forceinline__ int self_segment(const RunDev& R)
{
    if (RMAX == 1) return 13;
    const int side = 2 * R.reach + 1;
    return (side * side * side - 1) / 2;
}

template <bool MIXED, int RMAX>
__device__ __forceinline__ void build_nblist(const RunDev& R, int lcell, const uint32_t* __restrict__ cs,
                                             NbListR<RMAX>& L)
{
    constexpr int NS = RMAX == 1 ? 27 : 125;
    const int lane = threadIdx.x;
    const int c[3] = {lcell % R.n[0], (lcell / R.n[0]) % R.n[1], lcell / (R.n[0] * R.n[1])};
    const int side = RMAX == 1 ? 3 : 2 * R.reach + 1;
    const int nseg = side * side * side;
    int carry = 0;
#pragma unroll
    for (int base = 0; base < NS; base += 64) {
        const int seg = base + lane;
        int cnt = 0, st = 0;
        float sh[3] = {0.f, 0.f, 0.f};
        if (seg < nseg) {
            int cc;
            if (nb_cell<MIXED, RMAX>(R, c, seg, cc, sh)) {
                st = (int)cs[cc];
                cnt = (int)cs[cc + 1] - st;
            }
        }
        const int incl = wave_incl_scan(cnt) + carry;
        if (seg < NS) {
            L.start[seg] = st;
            L.off[seg] = seg < nseg ? incl - cnt : 0;  // past the stencil: overwritten with the total below
            L.sh[seg][0] = sh[0];
            L.sh[seg][1] = sh[1];
            L.sh[seg][2] = sh[2];
        }
        carry = __shfl(incl, 63, 64);
    }
    // the total after the last stencil entry, and in every entry past the run's stencil
    for (int k = nseg + lane; k <= NS; k += 64) L.off[k] = carry;
    __syncthreads();
}

// segment of virtual index v (largest k with off[k] <= v, over non-decreasing off[0..NS-1])
template <int NS>
__device__ __forceinline__ int nb_segment(const NbListT<NS>& L, int v)
{
    int k = 0;
#pragma unroll
    for (int step = NS > 64 ? 64 : 16; step >= 1; step >>= 1) {
        int t = k + step;
        if (t < NS && L.off[t] <= v) k = t;
    }
    return k;
}

// ------------------------------------------------------------------------------------------------
// Candidate pre-filter of the chunked pair passes. A candidate farther than the support radius from
// the bounding box of the supercell's targets is farther than that from every target, so the cull
// would reject it for all of them. Both passes drop such candidates before staging and stage the
// others in the same order, so the force pass replays the density pass's masks slot by slot as
// before. With supercells of edge ~H about half of the 27-cell stencil goes (MHD turbulence at 2M
// particles: 381 -> 194 candidates per supercell in a CPU replay), which halves the cull and the
// staging and leaves fewer, fuller chunks for the pair loops. The density pass filters; the force
// pass reads the kept list back from the mask row (recomputing it only where it was not stored),
// 11% faster than with no filter; the density pass, bound by latency more than by its cull, is 3%
// slower.
// ------------------------------------------------------------------------------------------------
constexpr int kPreCap = 768;  // larger stencils, and supercells of more than 64 particles, stage whole

struct PreList {
    uint16_t v[kPreCap];  // the kept candidates in list order: list index, and for the 27-cell stencil
                          // the segment in bits 10..14 (so staging needs no search)
    uint16_t self_s[64];  // kept index of each target of the supercell (a target is always kept)
};

// A kept-list entry: the list index (< kPreCap <= 1024) and, for the 27-cell stencil, its segment.
template <int NS>
__device__ __forceinline__ uint16_t pre_entry(int v, int sg)
{
    return (uint16_t)(NS == 27 ? v | (sg << 10) : v);
}

// List index and stencil segment of staged candidate su: from the kept list when the filter is on
// (the segment packed in the entry for the 27-cell stencil, else searched), else su itself.
template <int NS>
__device__ __forceinline__ void staged_candidate(const NbListT<NS>& L, const PreList& P, bool pre, int su, int& v,
                                                 int& sg)
{
    if (pre) {
        const int e = P.v[su];
        v = NS == 27 ? (e & 1023) : e;
        sg = NS == 27 ? (e >> 10) : nb_segment(L, v);
    } else {
        v = su;
        sg = nb_segment(L, v);
    }
}

__device__ __forceinline__ float wave_min(float x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = fminf(x, __shfl_xor(x, off, 64));
    return x;
}

__device__ __forceinline__ float wave_max(float x)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, 64));
    return x;
}

// Bounding box of the targets among a supercell's nt <= 64 particles (from their position records;
// empty, lo > hi, without targets). Its inputs need no neighbour list: a pass requests them first
// (target_box_load) and reduces them once the list is built (target_box), so the loads overlap.
struct TargetBox {
    float lx, ly, lz, hx, hy, hz;
};

struct TargetBoxInput {
    uint8_t species = 0, mask = 0;
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
};

__device__ __forceinline__ TargetBoxInput target_box_load(const StateBuf& S, const WorkBuf& W, int t0, int nt)
{
    const int lane = threadIdx.x;
    TargetBoxInput in;
    if (lane < nt) {
        const int i = t0 + lane;
        in.species = S.species[i];
        in.mask = S.mask[i];
        in.o = W.P0[i];  // a target is a source: its record holds its position
    }
    return in;
}

__device__ __forceinline__ TargetBox target_box(const TargetBoxInput& in, int nt)
{
    const bool has = (int)threadIdx.x < nt && is_target(in.species, in.mask);
    TargetBox b;
    b.lx = wave_min(has ? in.o.x : INFINITY);
    b.hx = wave_max(has ? in.o.x : -INFINITY);
    b.ly = wave_min(has ? in.o.y : INFINITY);
    b.hy = wave_max(has ? in.o.y : -INFINITY);
    b.lz = wave_min(has ? in.o.z : INFINITY);
    b.hz = wave_max(has ? in.o.z : -INFINITY);
    return b;
}

// Keeps in P the candidates of the supercell's list (of `total`) that lie within the support radius
// of its targets' bounding box B, and returns how many there are; with the filter off (`on` false)
// returns `total` and leaves P unused.
template <int NS>
__device__ __forceinline__ int prefilter(const NbListT<NS>& L, const WorkBuf& W, const TargetBox& B, int nt,
                                         int total, int self_v0, float H2, PreList& P, bool& on)
{
    on = nt <= 64 && total <= kPreCap;
    if (!on) return total;
    const int lane = threadIdx.x;
    const float lx = B.lx, hx = B.hx, ly = B.ly, hy = B.hy, lz = B.lz, hz = B.hz;
    // The test cannot drop a candidate the cull accepts, at any coordinate magnitude: per axis the
    // rounded distance to the box is at most the rounded distance to any target in it (rounding is
    // monotone), and both sums of squares round in the same fma order, so the cull's r2 is never
    // below this d2. The 2^-13 margin on H^2 is kept on top of that.
    const float H2m = H2 * (1.0f + 1.0f / 8192.0f);
    int count = 0;
    // four candidates per lane and round, all four records requested before any is tested
    for (int vb = 0; vb < total; vb += 256) {
        float4 q[4];
        int sg[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int v = vb + 64 * r + lane;
            sg[r] = 0;
            q[r] = make_float4(kSentinel, kSentinel, kSentinel, 0.f);
            if (v < total) {
                sg[r] = nb_segment(L, v);
                q[r] = W.P0[L.start[sg[r]] + (v - L.off[sg[r]])];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int v = vb + 64 * r + lane;
            bool keep = false;
            if (v < total) {
                const float px = q[r].x + L.sh[sg[r]][0], py = q[r].y + L.sh[sg[r]][1], pz = q[r].z + L.sh[sg[r]][2];
                const float dx = fmaxf(fmaxf(lx - px, px - hx), 0.0f);
                const float dy = fmaxf(fmaxf(ly - py, py - hy), 0.0f);
                const float dz = fmaxf(fmaxf(lz - pz, pz - hz), 0.0f);
                keep = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < H2m;  // sentinel records: +inf, dropped
            }
            const uint64_t b = __ballot(keep);
            if (keep) {
                const int sidx = count + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
                P.v[sidx] = pre_entry<NS>(v, sg[r]);
                const unsigned ts = (unsigned)(v - self_v0);
                if (ts < (unsigned)nt) P.self_s[ts] = (uint16_t)sidx;
            }
            count += __popcll(b);
        }
    }
    __syncthreads();
    return count;
}

// Cull lim (<= 32) consecutive lane candidates starting at LDS slot e0 (stride K): bit ii set iff
// r2 < H2, read off the sign of r2 - H2 (exact: the rounded difference has the sign of the exact
// one). Returns LSB-first bits.
__device__ __forceinline__ uint32_t cull32(const float4* __restrict__ sp, int e0, int K, int lim, float xi, float yi,
                                           float zi, float H2)
{
    uint32_t m = 0;
    if (lim == 32) {
        // candidates two at a time, both LDS reads issued before either is used: the scheduler
        // otherwise waits out every read alone to save registers (density pass 1.5% faster)
#pragma unroll
        for (int i0 = 0; i0 < 32; i0 += 2) {
            const float4 q0 = sp[e0 + i0 * K], q1 = sp[e0 + (i0 + 1) * K];
            __builtin_amdgcn_sched_barrier(0);
            float dx, dy, dz;
            const float r2a = geom_r2(xi, yi, zi, q0, dx, dy, dz);
            m = __builtin_amdgcn_alignbit(m, __float_as_uint(r2a - H2), 31);  // (m << 1) | sign
            const float r2b = geom_r2(xi, yi, zi, q1, dx, dy, dz);
            m = __builtin_amdgcn_alignbit(m, __float_as_uint(r2b - H2), 31);
        }
        return __builtin_bitreverse32(m);
    }
    for (int ii = 0; ii < lim; ++ii) {
        float dx, dy, dz;
        const float r2 = geom_r2(xi, yi, zi, sp[e0 + ii * K], dx, dy, dz);
        m = __builtin_amdgcn_alignbit(m, __float_as_uint(r2 - H2), 31);  // (m << 1) | sign
    }
    return __builtin_bitreverse32(m) >> (32 - lim);
}

// Staged order of a supercell's candidates in the chunked pair passes. The virtual list (the
// stencil's segments back to back) is cut into blocks of 8 consecutive candidates; the full blocks
// are laid out as 8 rows of consecutive blocks (the first `rem` rows one block longer) and staged
// column by column, the partial last block after them. A chunk of 64 staged slots then holds 8
// blocks from 8 places spread over the whole neighbourhood instead of a few neighbour supercells,
// so the accepted pairs of a chunk spread more evenly over the targets and the chunk's pair loop
// waits less for its busiest lane (a CPU replay of the force pass's lanes at 262,144 particles:
// 41.5 -> 34.0 pair rounds per wave), while the 8 lanes of a block still read 8 consecutive
// records (one candidate per slot, spread that way, made the staging loads uncoalesced and the
// force pass 14% slower). Both passes stage in this order, so the density pass's masks, indexed by
// staged slot, replay unchanged.
struct StageOrder {
    int nfull;  // full blocks of 8
    int cm1;    // blocks per short row - 1 (0 when nfull <= 8: the identity)
    int rem;    // rows of cm1 + 1 blocks

    __device__ __forceinline__ explicit StageOrder(int total)
    {
        nfull = total >> 3;
        const int c = (nfull + 7) >> 3;
        cm1 = c > 0 ? c - 1 : 0;
        rem = nfull - 8 * cm1;
    }
    // list index of staged slot u < total
    __device__ __forceinline__ int list_index(int u) const
    {
        if (u >= 8 * nfull) return u;
        const int b = u >> 3, i = b & 7, j = b >> 3;
        return 8 * (i * cm1 + (i < rem ? i : rem) + j) + (u & 7);
    }
    // staged slot of list index v < total
    __device__ __forceinline__ int slot(int v) const
    {
        if (v >= 8 * nfull) return v;
        const int blk = v >> 3, c = cm1 + 1;
        int i, j;
        if (blk < rem * c) {
            i = blk / c;
            j = blk - i * c;
        } else {  // only reached when cm1 > 0
            i = rem + (blk - rem * c) / cm1;
            j = blk - (i * cm1 + rem);
        }
        return 8 * (8 * j + i) + (v & 7);
    }
};

// Lane layout of a supercell-wave: nt targets, K = floor(64/nt) (capped) lanes per target.
struct LaneMap {
    int K, tpg, tl, k, gfirst, glast;
    bool inwave;
};

__device__ __forceinline__ LaneMap lane_map(int nt, int kmax)
{
    LaneMap m;
    const int lane = threadIdx.x;
    int K = nt >= 64 ? 1 : 64 / nt;
    m.K = K < kmax ? K : kmax;
    m.tpg = 64 / m.K;
    m.tl = (int)(((float)lane + 0.5f) * (1.0f / (float)m.K));  // lane / K, exact for lane < 64
    m.k = lane - m.tl * m.K;
    m.inwave = m.tl < m.tpg;
    m.gfirst = m.inwave ? m.tl * m.K : 63;
    m.glast = m.inwave ? m.gfirst + m.K - 1 : 63;
    return m;
}

// Whether a supercell-wave keeps its masks: one target group and every lane's candidates fit.
__device__ __forceinline__ bool masks_fit(int nt, int total, const LaneMap& m, int words = kMaskWords)
{
    return nt <= m.tpg && (total + m.K - 1) / m.K <= 32 * words;
}

// The pre-filter's kept list travels from the density pass to the force pass in the top of the
// supercell's mask row (kMaskWords x 64 words, seen as 16-bit entries): entry 1023 holds the kept
// count (or kNoList), entries 1022, 1021, ... the kept list indices. With the filter on, masks use at
// most kMaskWords - 1 words of each lane, so the count never shares a word with mask bits, and the
// list is stored only where it stays above the words the masks use.
constexpr uint16_t kNoList = 0xffffu;
constexpr int kRow16 = kMaskWords * 64 * 2;  // 16-bit entries per mask row

__device__ __forceinline__ bool list_fits(int count, int K)
{
    const int mask_words = ((count + K - 1) / K + 31) / 32;
    return count <= kRow16 - 1 - 128 * mask_words;
}

// Mask words are indexed by lane candidate c (candidate v = c*K + k of sub-lane k), so the two
// passes may stage with different chunk sizes. Bits for [c0, c0 + lim) may straddle two words.
// Words at or past `wlim` (those of candidates past the lane's last one: the last chunk's padding,
// all-zero bits) are not written, so the row above the lane's masks stays free.
struct MaskWriter {
    uint32_t* row = nullptr;
    uint32_t pend = 0;
    int pend_w = -1;
    int wlim = kMaskWords;  // <= kMaskWords
    __device__ __forceinline__ void put(uint32_t bits, int c0, int lim)
    {
        const int w = c0 >> 5, off = c0 & 31;
        if (w >= wlim) return;
        if (w != pend_w) {
            if (pend_w >= 0) row[pend_w * 64] = pend;
            pend = 0;
            pend_w = w;
        }
        pend |= bits << off;
        if (off + lim > 32) {
            row[w * 64] = pend;
            pend = bits >> (32 - off);
            pend_w = w + 1 < wlim ? w + 1 : -1;
        }
    }
    __device__ __forceinline__ void flush()
    {
        if (pend_w >= 0) row[pend_w * 64] = pend;
        pend_w = -1;
    }
};

__device__ __forceinline__ uint32_t mask_get(const uint32_t* row, int c0, int lim, int per_lane)
{
    int nvalid = per_lane - c0;
    if (nvalid <= 0) return 0u;
    if (nvalid > lim) nvalid = lim;
    const int w = c0 >> 5, off = c0 & 31;
    if (w >= kMaskWords) return 0u;
    uint32_t a = row[w * 64] >> off;
    if (off && off + nvalid > 32 && w + 1 < kMaskWords) a |= row[(w + 1) * 64] << (32 - off);
    return nvalid == 32 ? a : (a & ((1u << nvalid) - 1u));
}

// Compact the accepted LDS slots of each target's K lanes into lst[gfirst*32 ...) (in sub-lane,
// then candidate order) and return the target's count; the K lanes then share them round-robin.
template <class LT>
__device__ __forceinline__ int compact_group(uint32_t msk, int sb, const LaneMap& m, LT* lst)
{
    const int na = __popc(msk);
    const int incl = wave_incl_scan(na);
    const int bprev = __shfl(incl, m.gfirst > 0 ? m.gfirst - 1 : 0, 64);
    const int before = m.gfirst > 0 ? bprev : 0;
    const int gtot = __shfl(incl, m.glast, 64) - before;
    int wpos = m.gfirst * 32 + (incl - na - before);
    while (msk) {
        const int ii = __builtin_ctz(msk);
        msk &= msk - 1;
        lst[wpos++] = (LT)(m.k + (sb + ii) * m.K);
    }
    return m.inwave ? gtot : 0;
}

// K partial sums of a target combined in lane order (deterministic); then one lane per particle
// writes rho, the neighbour count and the interaction records of the force pass, and the run's
// fast-speed maximum (cleaning speed) is raised.
template <bool MHD>
__device__ __forceinline__ void density_finish(const RunDev& R, int r, StateBuf& S, WorkBuf& W, const LaneMap& lm,
                                               float acc, uint32_t cnt, bool valid, bool tgt, uint8_t spc,
                                               uint8_t live, int i, float m_own, float h, float invh, float uu,
                                               float vx, float vy, float vz, float rho_in, float bx, float by,
                                               float bz, float psi, float gam, int kern)
{
    const int K = lm.K, k = lm.k;
    const int lane = threadIdx.x;
    // combine the K partial sums in lane order (deterministic)
    float acc_t = 0.0f;
    uint32_t cnt_t = 0;
    for (int kk = 0; kk < K; ++kk) {
        const int src = lm.gfirst + kk < 64 ? lm.gfirst + kk : 63;
        acc_t += __shfl(acc, src, 64);
        cnt_t += __shfl(cnt, src, 64);
    }
    // epilogue: one lane per particle writes rho and its interaction records
    float cmax = 0.0f;
    if (k == 0 && valid && is_source(spc, live)) {
#pragma clang fp contract(off)
        const float sig = R.sigma;
        const float wfac = sig / (h * h * h);
        const float gfac = (kern == 2) ? -5.0f * sig / (h * h * h * h * h) : sig / (h * h * h * h);
        float rho;
        if (tgt) {
            rho = wfac * (acc_t + m_own * R.f0);
            S.rho[i] = rho;
            W.nnb[i] = cnt_t;
        } else {
            rho = rho_in;
        }
        const float P = (gam - 1.0f) * rho * uu;
        const float invrho = 1.0f / rho;
        const float B2 = bx * bx + by * by + bz * bz;
        const float bdr = invrho * invrho;
        const float ph = P * bdr;
        const float pr = (P + 0.5f * B2) * bdr;
        const float psr = psi * bdr;
        const float c = sqrtf((gam * P + B2) * invrho);
        W.P1[i] = make_float4(vx, vy, vz, c);
        W.P2[i] = make_float4(h, invh, gfac, rho);
        W.P3[i] = make_float4(pr, ph, bdr, psr);
        if (MHD) W.P4[i] = make_float4(bx, by, bz, psi);
        if (tgt) cmax = c;
    }
    if (k == 0 && valid && !tgt) W.nnb[i] = 0;
    if (MHD && R.mhd && R.cleaning) {
        for (int off = 32; off >= 1; off >>= 1) cmax = fmaxf(cmax, __shfl_xor(cmax, off, 64));
        // every wave of the run targets this one word: test before the atomic, so that only a
        // wave holding a new maximum pays for one (the value only grows, so a stale read costs
        // at most a redundant atomic)
        if (lane == 0 && cmax > 0.0f) {
            const uint32_t bits = __float_as_uint(cmax);
            if (bits > __atomic_load_n(&W.chbits[r], __ATOMIC_RELAXED)) atomicMax(&W.chbits[r], bits);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// density pass: cull of the 27-supercell candidates staged in LDS, density sum, candidate masks,
// and the per-particle interaction records of the force pass
// ------------------------------------------------------------------------------------------------
// With the Wendland kernel the pass fits 64 VGPRs without spilling: 8 waves per SIMD instead of 7,
// 5% faster on the 2M-particle turbulence step. The other kernels keep the compiler's choice
// (a cap spills them), as do the balanced pass and the 125-cell stencil with 256-candidate chunks,
// whose LDS allows 6.
template <int CH, bool MIXED, bool MHD, int KT, int RMAX>
__device__ __forceinline__ void density_wave(int vb, const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                             int nblk_total, StateBuf& S, WorkBuf& W, int kmax, NbListR<RMAX>& L,
                                             float4 (&sp)[CH], PreList& P)
{
    constexpr int NS = RMAX == 1 ? 27 : 125;
    const int cell = xcd_cell(vb, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    // the targets' box inputs are requested before the neighbour list, so their loads overlap the list's
    const TargetBoxInput box_in = nt <= 64 ? target_box_load(S, W, t0, nt) : TargetBoxInput{};
    build_nblist<MIXED, RMAX>(R, cell - R.c0, W.cs, L);
    const int self_v0 = L.off[self_segment<RMAX>(R)];
    bool pre;
    const int total = prefilter<NS>(L, W, nt <= 64 ? target_box(box_in, nt) : TargetBox{}, nt, L.off[NS], self_v0,
                                    R.H2, P, pre);  // staged candidates
    const StageOrder so(total);
    const int lane = threadIdx.x;
    const LaneMap lm = lane_map(nt, kmax);
    const int K = lm.K, k = lm.k;
    const int per = CH / K;   // lane candidates per chunk
    const int chk = per * K;  // staged entries per chunk (a multiple of K)
    const int kern = KT >= 0 ? KT : R.kernel;  // compile-time when every run uses one kernel
    const float H2 = R.H2;
    const float gam = R.gamma;
    // no mask table: density only; with the filter on the top word of each lane carries the kept list
    const bool keep = masks_fit(nt, total, lm, pre ? kMaskWords - 1 : kMaskWords) && W.nbmask != nullptr;
    MaskWriter mw;
    mw.row = W.nbmask + (size_t)cell * (kMaskWords * 64) + lane;
    mw.pend = 0;
    mw.pend_w = -1;
    mw.wlim = ((total + K - 1) / K + 31) / 32;  // the words the lane's candidates need

    for (int tg = 0; tg < nt; tg += lm.tpg) {
        const int t = tg + lm.tl;
        const bool valid = lm.inwave && t < nt;
        const int i = t0 + (valid ? t : 0);
        const uint8_t live = S.mask[i], spc = S.species[i];
        const bool tgt = valid && is_target(spc, live);
        const float4 own = W.P0[i];
        const float h = S.h[i];
        // inputs of the epilogue, loaded early so their latency overlaps the cull
        const float uu = S.u[i];
        const float vx = S.vel[0][i], vy = S.vel[1][i], vz = S.vel[2][i];
        const float rho_in = S.rho[i];
        float bx = 0.f, by = 0.f, bz = 0.f, psi = 0.f;
        if (MHD && R.mhd) {
            bx = S.B[0][i];
            by = S.B[1][i];
            bz = S.B[2][i];
            psi = S.psi[i];
        }
        const float invh = 1.0f / h;
        const int self_u = tgt ? so.slot(pre ? (int)P.self_s[t] : self_v0 + t) : -1;  // staged slot of the self pair
        float acc = 0.0f;
        uint32_t cnt = 0;

        for (int base = 0; base < total; base += chk) {
            for (int e = lane; e < chk; e += 64) {
                const int u = base + e;
                float4 p = make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
                if (u < total) {
                    int v, sg;
                    staged_candidate<NS>(L, P, pre, so.list_index(u), v, sg);
                    p = W.P0[L.start[sg] + (v - L.off[sg])];
                    p.x += L.sh[sg][0];
                    p.y += L.sh[sg][1];
                    p.z += L.sh[sg][2];
                }
                sp[e] = p;
            }
            __syncthreads();
            const int cbase = base / K;
            // lane slots of this chunk that hold candidates (the most any lane has): the last chunk's
            // padding is not culled
            const int nval = ((total - base < chk ? total - base : chk) + K - 1) / K;
            for (int sb = 0; sb < nval; sb += 32) {
                const int lim = nval - sb < 32 ? nval - sb : 32;
                uint32_t msk = cull32(sp, k + sb * K, K, lim, own.x, own.y, own.z, H2);
                // the self pair (r2 = 0) passes the cull: drop it here; its density term is the
                // self term m_i W(0, h_i) added below, as DensityInitSelf does
                const int sc = self_u - k;
                if (sc >= 0 && sc % K == 0) {
                    const int bit = sc / K - cbase - sb;
                    if (bit >= 0 && bit < lim) msk &= ~(1u << bit);
                }
                if (!tgt) msk = 0;
                if (keep) mw.put(msk, cbase + sb, lim);
                cnt += (uint32_t)__popc(msk);
                // each lane sums its own accepted candidates (the density pairs are too cheap for
                // sharing them among a target's lanes to pay for the compaction)
                const int kb = (int)sizeof(float4) * K;
                const int eb0 = (int)sizeof(float4) * (k + sb * K);  // byte offset of bit 0's slot
                while (msk) {
                    const int ii = __builtin_ctz(msk);
                    msk &= msk - 1;
                    const float4 p = lds_at(sp, (int)__umul24((unsigned)ii, (unsigned)kb) + eb0);
                    float dx, dy, dz;
                    const float r2 = geom_r2(own.x, own.y, own.z, p, dx, dy, dz);
                    acc += p.w * kshape(kern, fast_sqrt(r2) * invh);
                }
            }
            __syncthreads();
        }
        if (keep) mw.flush();
        density_finish<MHD>(R, r, S, W, lm, acc, cnt, valid, tgt, spc, live, i, own.w, h, invh, uu, vx, vy, vz,
                            rho_in, bx, by, bz, psi, gam, kern);
    }
    // the kept list for the force pass, stored last: stores count in the same wait counter as the
    // loads, so stored earlier they would delay every later wait for the staging loads
    if (pre && W.nbmask != nullptr) {
        uint16_t* row16 = reinterpret_cast<uint16_t*>(W.nbmask + (size_t)cell * (kMaskWords * 64));
        const bool store = keep && list_fits(total, K);
        if (store)
            for (int e = lane; e < total; e += 64) row16[kRow16 - 2 - e] = P.v[e];
        if (lane == 0) row16[kRow16 - 1] = store ? (uint16_t)total : kNoList;
    }
}

template <int CH, bool MIXED, bool MHD, int KT, int RMAX = 1, bool GRID = false>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu((KT == 2 && !(RMAX == 2 && CH > 128)) ? 8 : 1)))  // KT 2: Wendland
void k_density(const RunDev* __restrict__ runs, int nruns, int ncell_tot, int nblk_total, StateBuf S, WorkBuf W,
               int kmax)
{
    __shared__ NbListR<RMAX> L;
    __shared__ float4 sp[CH];
    __shared__ PreList P;
    if constexpr (GRID) {
        // a persistent grid: each wave walks the supercells at a stride of the grid (a multiple
        // of 8, so every cell a wave visits belongs to its own XCD's range)
        for (int vb = blockIdx.x; vb < nblk_total; vb += gridDim.x)
            density_wave<CH, MIXED, MHD, KT, RMAX>(vb, runs, nruns, ncell_tot, nblk_total, S, W, kmax, L, sp, P);
    } else {
        // one supercell per wave
        density_wave<CH, MIXED, MHD, KT, RMAX>((int)blockIdx.x, runs, nruns, ncell_tot, nblk_total, S, W, kmax, L,
                                               sp, P);
    }
}

// ------------------------------------------------------------------------------------------------
// force pass: pressure gradient, P dV, Monaghan viscosity, SPMHD stress, induction, cleaning
// ------------------------------------------------------------------------------------------------
struct ForceAcc {
    float ax, ay, az, DV, AVU, vmax;
    float Ti, Tj, TBx, TBy, TBz, VBx, VBy, VBz, DB, PSx, PSy, PSz;
};

// ONE_H: every particle of the pair's run has the same smoothing length h; the pair then takes
// s_j = s_i and the mean h = h without testing or forming them (the same bits: 0.5 (h + h) = h).
template <bool MHD, bool ONE_H = false>
__device__ __forceinline__ void force_pair(const float4& o0, const float4& o1, const float4& o2, const float4& o3,
                                           const float4& o4, const float4& q0, const float4& q1, const float4& q2,
                                           const float4& q3, const float4& q4, int kern, float alpha, float beta,
                                           bool mhd, bool clean, ForceAcc& A)
{
    float dx, dy, dz;
    const float r2 = geom_r2(o0.x, o0.y, o0.z, q0, dx, dy, dz);
    const float invR = r2 > 0.0f ? fast_rsq(r2) : 0.0f;
    const float rr = r2 * invR;
    const float s_i = kgrad(kern, rr, invR, rr * o2.y, o2.z);
    // equal smoothing lengths (the records hold h, 1/h and the gradient factor derived from it)
    // give the same gradient bit for bit: the second evaluation is skipped then
    float s_j = s_i;
    if (!ONE_H && __builtin_expect(q2.x != o2.x, 0)) s_j = kgrad(kern, rr, invR, rr * q2.y, q2.z);
    const float dvx = o1.x - q1.x, dvy = o1.y - q1.y, dvz = o1.z - q1.z;
    const float vr = dvx * dx + dvy * dy + dvz * dz;
    const float m = q0.w;
    const float ms_i = m * s_i, ms_j = m * s_j;
    const float fiso = o3.x * ms_i + q3.x * ms_j;
    A.ax -= fiso * dx;
    A.ay -= fiso * dy;
    A.az -= fiso * dz;
    A.DV += ms_i * vr;
    const float w = vr * invR;
    const float vs = o1.w + q1.w - beta * fminf(w, 0.0f);
    // signal speeds are >= +0, so their float bits order as integers: an integer max is the float
    // max without the NaN-quieting copy fmaxf needs for its loop-carried operand
    A.vmax = __int_as_float(max(__float_as_int(A.vmax), __float_as_int(vs)));
    if (alpha > 0.0f) {  // uniform per run; receding pairs get mu = 0, hence Pi = 0
        const float hb = ONE_H ? o2.x : 0.5f * (o2.x + q2.x);
        const float mu = hb * fminf(vr, 0.0f) * fast_rcp(r2 + 0.01f * hb * hb);
        // Pi = (beta mu - alpha cb) mu / rb, cb and rb the pair's mean sound speed and density, and
        // gv = Pi (m s_i + m s_j) / 2. The halves are power-of-two scalings, exact in FP32: alpha cb
        // = (alpha / 2)(c_i + c_j), and (1 / rb) / 2 = 1 / (rho_i + rho_j), so Ph = Pi / 2.
        const float Ph = (beta * mu - (0.5f * alpha) * (o1.w + q1.w)) * mu * fast_rcp(o2.w + q2.w);
        const float gv = Ph * (ms_i + ms_j);
        A.ax -= gv * dx;
        A.ay -= gv * dy;
        A.az -= gv * dz;
        A.AVU += gv * vr;
    }
    if (MHD && mhd) {
        const float Bir = o4.x * dx + o4.y * dy + o4.z * dz;
        const float Bjr = q4.x * dx + q4.y * dy + q4.z * dz;
        const float ti = Bir * o3.z * ms_i;
        const float tj = Bjr * q3.z * ms_j;
        A.Ti += ti;
        A.Tj += tj;
        A.TBx += tj * q4.x;
        A.TBy += tj * q4.y;
        A.TBz += tj * q4.z;
        const float wv = ms_i * Bir;
        A.VBx += wv * dvx;
        A.VBy += wv * dvy;
        A.VBz += wv * dvz;
        A.DB += ms_i * (Bir - Bjr);
        if (clean) {
            const float fps = o3.w * ms_i + q3.w * ms_j;
            A.PSx += fps * dx;
            A.PSy += fps * dy;
            A.PSz += fps * dz;
        }
    }
}

// One target's derivatives, signal speed and div B from its summed pair terms T.
template <bool MHD>
__device__ __forceinline__ void force_write(const RunDev& R, WorkBuf& W, const ForceAcc& T, int i, const float4& o1,
                                            const float4& o2, const float4& o3, const float4& o4, bool mhd,
                                            bool clean, float borve, float ch)
{
    const float rho = o2.w;
    const float invrho = 1.0f / rho;
    float ax = T.ax, ay = T.ay, az = T.az;
    if (MHD && mhd) {
        const float tcoef = T.Ti - borve * (T.Ti + T.Tj);
        ax = ax + T.TBx + o4.x * tcoef;
        ay = ay + T.TBy + o4.y * tcoef;
        az = az + T.TBz + o4.z * tcoef;
    }
    W.dvdt[0][i] = ax;
    W.dvdt[1][i] = ay;
    W.dvdt[2][i] = az;
    W.dudt[i] = o3.y * T.DV + 0.5f * T.AVU;
    W.vsig[i] = fmaxf(2.0f * o1.w, T.vmax);
    if (MHD && mhd) {
        const float divB = -T.DB * invrho;
        const float divv = -T.DV * invrho;
        float dbx = -(T.VBx - o4.x * T.DV) * invrho;
        float dby = -(T.VBy - o4.y * T.DV) * invrho;
        float dbz = -(T.VBz - o4.z * T.DV) * invrho;
        float dpsi = 0.0f;
        if (clean) {
            dbx -= rho * T.PSx;
            dby -= rho * T.PSy;
            dbz -= rho * T.PSz;
            const float psi = o4.w;
            dpsi = -ch * ch * divB - psi * R.clean_sigma * ch * o2.y - 0.5f * psi * divv;
        }
        W.dbdt[0][i] = dbx;
        W.dbdt[1][i] = dby;
        W.dbdt[2][i] = dbz;
        W.dpsidt[i] = dpsi;
        W.divb[i] = divB;
    }
}

// x from the lane D below in the same 16-lane row (DPP row_shr:D; lanes at the row's start read 0)
template <int D>
__device__ __forceinline__ float row_shr(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x110 + D, 0xf, 0xf, true));
}

// K partial force sums of a target combined (deterministic), field by field; then one lane per
// target writes its derivatives, signal speed and div B. For K = 2, 4, 8 or 16 a target's lanes lie
// in one 16-lane row, and log2 K DPP steps (each lane adds the partial of the lane 1, 2, 4, ... below
// it) leave the target's total in its last lane; other K take the shuffles of each lane's partial
// in lane order to every lane of the target.
template <bool MHD>
__device__ __forceinline__ void force_finish(const RunDev& R, WorkBuf& W, const LaneMap& lm, ForceAcc& A, bool tgt,
                                             int i, const float4& o1, const float4& o2, const float4& o3,
                                             const float4& o4, bool mhd, bool clean, float borve, float ch)
{
    const int K = lm.K, k = lm.k;
    constexpr int NF = MHD ? 17 : 5;
    float v[NF] = {A.ax, A.ay, A.az, A.DV, A.AVU};
    if constexpr (MHD) {
        const float m[12] = {A.Ti, A.Tj, A.TBx, A.TBy, A.TBz, A.VBx, A.VBy, A.VBz, A.DB, A.PSx, A.PSy, A.PSz};
#pragma unroll
        for (int f = 0; f < 12; ++f) v[5 + f] = m[f];
    }
    float s[NF];
    float vm;
    bool writer;
    if (K == 1 || K == 2 || K == 4 || K == 8 || K == 16) {
        auto step = [&](auto d) {
            constexpr int D = decltype(d)::value;
#pragma unroll
            for (int f = 0; f < NF; ++f) v[f] += row_shr<D>(v[f]);
            // signal speeds are >= +0: an integer max on their bits
            A.vmax = __int_as_float(max(__float_as_int(A.vmax), __float_as_int(row_shr<D>(A.vmax))));
        };
        if (K >= 2) step(std::integral_constant<int, 1>{});
        if (K >= 4) step(std::integral_constant<int, 2>{});
        if (K >= 8) step(std::integral_constant<int, 4>{});
        if (K >= 16) step(std::integral_constant<int, 8>{});
#pragma unroll
        for (int f = 0; f < NF; ++f) s[f] = v[f];
        vm = A.vmax;
        writer = k == K - 1;
    } else {
        // each lane's partials shuffled to every lane of the target, added in lane order. All fields
        // of one lane step are fetched together, so a step's shuffles overlap instead of each waiting
        // for the last.
#pragma unroll
        for (int f = 0; f < NF; ++f) s[f] = 0.0f;
        vm = 0.0f;
        for (int kk = 0; kk < K; ++kk) {
            const int src = lm.gfirst + kk < 64 ? lm.gfirst + kk : 63;
            float t[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) t[f] = __shfl(v[f], src, 64);
            const float tm = __shfl(A.vmax, src, 64);
#pragma unroll
            for (int f = 0; f < NF; ++f) s[f] += t[f];
            vm = fmaxf(vm, tm);
        }
        writer = k == 0;
    }
    ForceAcc& T = A;
    T.ax = s[0];
    T.ay = s[1];
    T.az = s[2];
    T.DV = s[3];
    T.AVU = s[4];
    T.vmax = vm;
    if constexpr (MHD) {
        T.Ti = s[5];
        T.Tj = s[6];
        T.TBx = s[7];
        T.TBy = s[8];
        T.TBz = s[9];
        T.VBx = s[10];
        T.VBy = s[11];
        T.VBz = s[12];
        T.DB = s[13];
        T.PSx = s[14];
        T.PSy = s[15];
        T.PSz = s[16];
    }
    if (writer && tgt) force_write<MHD>(R, W, T, i, o1, o2, o3, o4, mhd, clean, borve, ch);
}

template <int CH, bool MIXED, bool MHD, int KT, int BAL, int RMAX, bool ONE_H = false>
__device__ __forceinline__ void forces_wave(int vb, const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                            int nblk_total, StateBuf& S, WorkBuf& W, int kmax, NbListR<RMAX>& L,
                                            float4 (&sp)[MHD ? 5 : 4][CH], ListT<CH> (&lst)[BAL == 1 ? 64 * 32 : 1],
                                            PreList& P)
{
    constexpr int NS = RMAX == 1 ? 27 : 125;
    const int cell = xcd_cell(vb, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    const int lane = threadIdx.x;
    // the density pass's kept set, stored in the mask row: its count and first 256 entries are
    // requested before the neighbour list is built, so their loads overlap its cell-start loads
    // (they are used only if this step's filter applies to the supercell, which the list decides)
    const uint16_t* row16 =
        W.nbmask != nullptr ? reinterpret_cast<const uint16_t*>(W.nbmask + (size_t)cell * (kMaskWords * 64)) : nullptr;
    int n_row = kNoList;
    uint16_t head[4] = {0, 0, 0, 0};
    if (row16 != nullptr && nt <= 64) {
        n_row = row16[kRow16 - 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) head[q] = row16[kRow16 - 2 - (lane + 64 * q)];
    }
    build_nblist<MIXED, RMAX>(R, cell - R.c0, W.cs, L);
    const int self_v0 = L.off[self_segment<RMAX>(R)];
    const LaneMap lm = lane_map(nt, kmax);
    const int K = lm.K, k = lm.k;
    // read back where the density pass stored it, else recomputed
    bool pre = nt <= 64 && L.off[NS] <= kPreCap;
    bool stored = false;
    int total = L.off[NS];
    if (pre && row16 != nullptr && n_row != kNoList) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (lane + 64 * q < n_row) P.v[lane + 64 * q] = head[q];
        for (int e = lane + 256; e < n_row; e += 64) P.v[e] = row16[kRow16 - 2 - e];
        __syncthreads();
        total = n_row;
        stored = true;
    }
    if (!stored)
        total = prefilter<NS>(L, W, pre ? target_box(target_box_load(S, W, t0, nt), nt) : TargetBox{}, nt, total,
                              self_v0, R.H2, P, pre);
    const StageOrder so(total);  // the density pass's staged order: its masks replay slot by slot
    const int per = CH / K;
    const int chk = per * K;
    const int kern = KT >= 0 ? KT : R.kernel;  // compile-time when every run uses one kernel
    const float H2 = R.H2;
    const float alpha = R.alpha, beta = R.beta;
    const bool mhd = MHD && R.mhd;
    const bool clean = mhd && R.cleaning;
    const float borve = mhd ? R.borve : 0.0f;
    const float ch = __uint_as_float(W.chbits[r]);
    const bool keep = masks_fit(nt, total, lm, pre ? kMaskWords - 1 : kMaskWords);
    const int per_lane = (total + K - 1) / K;
    const uint32_t* mrow = W.nbmask + (size_t)cell * (kMaskWords * 64) + lane;

    for (int tg = 0; tg < nt; tg += lm.tpg) {
        const int t = tg + lm.tl;
        const bool valid = lm.inwave && t < nt;
        const int i = t0 + (valid ? t : 0);
        const bool tgt = valid && is_target(S.species[i], S.mask[i]);
        const float4 o0 = W.P0[i], o1 = W.P1[i], o2 = W.P2[i], o3 = W.P3[i];
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MHD) o4 = W.P4[i];
        // the self slot is needed only where the pass culls (a stored list implies stored masks)
        const int self_u = tgt && !keep ? so.slot(pre ? (int)P.self_s[t] : self_v0 + t) : -1;
        ForceAcc A = {};

        for (int base = 0; base < total; base += chk) {
            for (int e = lane; e < chk; e += 64) {
                const int u = base + e;
                if (u < total) {
                    int v, sg;
                    staged_candidate<NS>(L, P, pre, so.list_index(u), v, sg);
                    const int j = L.start[sg] + (v - L.off[sg]);
                    float4 p = W.P0[j];
                    p.x += L.sh[sg][0];
                    p.y += L.sh[sg][1];
                    p.z += L.sh[sg][2];
                    sp[0][e] = p;
                    sp[1][e] = W.P1[j];
                    sp[2][e] = W.P2[j];
                    sp[3][e] = W.P3[j];
                    if constexpr (MHD) sp[4][e] = W.P4[j];
                } else {
                    sp[0][e] = make_float4(kSentinel, kSentinel, kSentinel, 0.0f);
                }
            }
            __syncthreads();
            const int cbase = base / K;
            for (int sb = 0; sb < per; sb += 32) {
                const int lim = per - sb < 32 ? per - sb : 32;
                uint32_t msk;
                if (keep) {
                    msk = mask_get(mrow, cbase + sb, lim, per_lane);
                } else {
                    msk = cull32(sp[0], k + sb * K, K, lim, o0.x, o0.y, o0.z, H2);
                    const int sc = self_u - k;  // the self pair adds exact zeros; drop it anyway
                    if (sc >= 0 && sc % K == 0) {
                        const int bit = sc / K - cbase - sb;
                        if (bit >= 0 && bit < lim) msk &= ~(1u << bit);
                    }
                }
                if (!tgt) msk = 0;
                if (BAL == 0) {
                    const int kb = (int)sizeof(float4) * K;
                    const int eb0 = (int)sizeof(float4) * (k + sb * K);  // byte offset of bit 0's slot
                    while (msk) {
                        const int ii = __builtin_ctz(msk);
                        msk &= msk - 1;
                        const int eb = (int)__umul24((unsigned)ii, (unsigned)kb) + eb0;
                        float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                        if constexpr (MHD) q4 = lds_at(sp[4], eb);
                        // all of the pair's records read before any is used: one LDS wait per pair
                        // instead of one per record (force pass 3% faster; same arithmetic)
                        const float4 q0 = lds_at(sp[0], eb), q1 = lds_at(sp[1], eb), q2 = lds_at(sp[2], eb),
                                     q3 = lds_at(sp[3], eb);
                        __builtin_amdgcn_sched_barrier(0);
                        force_pair<MHD, ONE_H>(o0, o1, o2, o3, o4, q0, q1, q2, q3, q4, kern, alpha, beta, mhd, clean,
                                               A);
                    }
                } else {
                    const int gtot = compact_group(msk, sb, lm, lst);
                    __syncthreads();
                    const ListT<CH>* mine = lst + lm.gfirst * 32;  // this target's accepted slots
                    for (int q = k; q < gtot; q += K) {
                        const int eb = (int)sizeof(float4) * (int)mine[q];
                        float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                        if constexpr (MHD) q4 = lds_at(sp[4], eb);
                        const float4 q0 = lds_at(sp[0], eb), q1 = lds_at(sp[1], eb), q2 = lds_at(sp[2], eb),
                                     q3 = lds_at(sp[3], eb);
                        __builtin_amdgcn_sched_barrier(0);
                        force_pair<MHD, ONE_H>(o0, o1, o2, o3, o4, q0, q1, q2, q3, q4, kern, alpha, beta, mhd, clean,
                                               A);
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
        }
        force_finish<MHD>(R, W, lm, A, tgt, i, o1, o2, o3, o4, mhd, clean, borve, ch);
    }
}

template <int CH, bool MIXED, bool MHD, int KT, int BAL, int RMAX = 1, bool GRID = false, bool ONE_H = false>
__global__ __launch_bounds__(64)
// the balanced loop fits 96 VGPRs without spilling (5 waves per SIMD; the compiler's own choice is 97)
__attribute__((amdgpu_waves_per_eu(BAL == 1 && !GRID ? 5 : 1)))
void k_forces(const RunDev* __restrict__ runs, int nruns, int ncell_tot, int nblk_total, StateBuf S, WorkBuf W,
              int kmax, int balance)
{
    constexpr int NP = MHD ? 5 : 4;
    __shared__ NbListR<RMAX> L;
    __shared__ float4 sp[NP][CH];
    __shared__ ListT<CH> lst[BAL == 1 ? 64 * 32 : 1];  // compaction lists of the balanced pair loop
    __shared__ PreList P;
    if constexpr (GRID) {
        // a persistent grid: each wave walks the supercells at a stride of the grid (a multiple
        // of 8, so every cell a wave visits belongs to its own XCD's range)
        for (int vb = blockIdx.x; vb < nblk_total; vb += gridDim.x)
            forces_wave<CH, MIXED, MHD, KT, BAL, RMAX>(vb, runs, nruns, ncell_tot, nblk_total, S, W, kmax, L, sp,
                                                       lst, P);
    } else {
        // one supercell per wave
        forces_wave<CH, MIXED, MHD, KT, BAL, RMAX, ONE_H>((int)blockIdx.x, runs, nruns, ncell_tot, nblk_total, S, W,
                                                          kmax, L, sp, lst, P);
    }
}

// ------------------------------------------------------------------------------------------------
// Whole-cell pair passes (tuning pair_balance = 2).
//
// Lane (target t, sub-lane k) owns the contiguous candidate range [k*Lr, min((k+1)*Lr, total)) of
// the 27-supercell virtual list, Lr = ceil(total / K). The density pass stages the positions of
// all candidates at once (chunks of CAPD only for very large stencils), culls its range and stores
// the accepted bits of the range (bit b <-> candidate k*Lr + b). Because the ranges of a target's
// K lanes are consecutive, concatenating their accepted candidates gives each target's neighbour
// list sorted by candidate index. The force pass builds these lists once per supercell in LDS, then
// stages the records chunk by chunk: within a chunk a target's pairs are one contiguous run of its
// list, dealt round-robin to its K lanes, with no per-chunk compaction. Cells whose lists do not
// fit (or whose masks were not stored) take a direct per-lane loop over their own bits.
// ------------------------------------------------------------------------------------------------
constexpr int kCellListCap = 2048;  // uint16 entries of the per-supercell neighbour lists

// Popcount of bits [lo, hi) of a lane's mask row (word w at row[w * 64]).
__device__ __forceinline__ int mask_count(const uint32_t* row, int lo, int hi)
{
    int n = 0;
    for (int b = lo; b < hi;) {
        const int w = b >> 5, off = b & 31;
        const int take = (32 - off) < (hi - b) ? (32 - off) : (hi - b);
        uint32_t m = row[w * 64] >> off;
        if (take < 32) m &= (1u << take) - 1u;
        n += __popc(m);
        b += take;
    }
    return n;
}

// Bits [lo, lo + lim) of a lane's mask row, LSB first (lim <= 32).
__device__ __forceinline__ uint32_t mask_bits(const uint32_t* row, int lo, int lim)
{
    const int w = lo >> 5, off = lo & 31;
    uint32_t a = row[w * 64] >> off;
    if (off && off + lim > 32) a |= row[(w + 1) * 64] << (32 - off);
    return lim == 32 ? a : (a & ((1u << lim) - 1u));
}

template <int CAPD, bool MIXED, bool MHD, int KT>
__global__ __launch_bounds__(64) void k_density_cell(const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                                     int nblk_total, StateBuf S, WorkBuf W, int kmax)
{
    constexpr int RMAX = 1, NS = 27;
    __shared__ NbListR<RMAX> L;
    __shared__ float4 sp[CAPD];
    const int cell = xcd_cell(blockIdx.x, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    build_nblist<MIXED, RMAX>(R, cell - R.c0, W.cs, L);
    const int total = L.off[NS];
    const int lane = threadIdx.x;
    const LaneMap lm = lane_map(nt, kmax);
    const int K = lm.K, k = lm.k;
    const int Lr = (total + K - 1) / K;
    const bool keep = nt <= lm.tpg && Lr <= 32 * kMaskWords && W.nbmask != nullptr;
    const int kern = KT >= 0 ? KT : R.kernel;  // compile-time when every run uses one kernel
    const float H2 = R.H2;
    const float gam = R.gamma;
    MaskWriter mw;
    mw.row = W.nbmask + (size_t)cell * (kMaskWords * 64) + lane;
    mw.pend = 0;
    mw.pend_w = -1;
    mw.wlim = kMaskWords;

    for (int tg = 0; tg < nt; tg += lm.tpg) {
        const int t = tg + lm.tl;
        const bool valid = lm.inwave && t < nt;
        const int i = t0 + (valid ? t : 0);
        const uint8_t live = S.mask[i], spc = S.species[i];
        const bool tgt = valid && is_target(spc, live);
        const float4 own = W.P0[i];
        const float h = S.h[i];
        const float uu = S.u[i];
        const float vx = S.vel[0][i], vy = S.vel[1][i], vz = S.vel[2][i];
        const float rho_in = S.rho[i];
        float bx = 0.f, by = 0.f, bz = 0.f, psi = 0.f;
        if (MHD && R.mhd) {
            bx = S.B[0][i];
            by = S.B[1][i];
            bz = S.B[2][i];
            psi = S.psi[i];
        }
        const float invh = 1.0f / h;
        const int self_v = L.off[self_segment<RMAX>(R)] + t;
        const int v0 = lm.inwave ? k * Lr : total;
        const int v1 = v0 + Lr < total ? v0 + Lr : total;
        float acc = 0.0f;
        uint32_t cnt = 0;
        for (int base = 0; base < total; base += CAPD) {
            const int top = base + CAPD < total ? base + CAPD : total;
            for (int v = base + lane; v < top; v += 64) {
                const int sg = nb_segment(L, v);
                float4 p = W.P0[L.start[sg] + (v - L.off[sg])];
                p.x += L.sh[sg][0];
                p.y += L.sh[sg][1];
                p.z += L.sh[sg][2];
                sp[v - base] = p;
            }
            __syncthreads();
            const int a = v0 > base ? v0 : base;
            const int b = v1 < top ? v1 : top;
            for (int cc = a; cc < b; cc += 32) {
                const int lim = b - cc < 32 ? b - cc : 32;
                uint32_t msk = cull32(sp, cc - base, 1, lim, own.x, own.y, own.z, H2);
                // the self pair passes the cull; its density term is the self term added at the end
                if (self_v >= cc && self_v < cc + lim) msk &= ~(1u << (self_v - cc));
                if (!tgt) msk = 0;
                if (keep) mw.put(msk, cc - v0, lim);
                cnt += (uint32_t)__popc(msk);
                while (msk) {
                    const int ii = __builtin_ctz(msk);
                    msk &= msk - 1;
                    const float4 p = sp[cc - base + ii];
                    float dx, dy, dz;
                    const float r2 = geom_r2(own.x, own.y, own.z, p, dx, dy, dz);
                    acc += p.w * kshape(kern, fast_sqrt(r2) * invh);
                }
            }
            __syncthreads();
        }
        if (keep) mw.flush();
        density_finish<MHD>(R, r, S, W, lm, acc, cnt, valid, tgt, spc, live, i, own.w, h, invh, uu, vx, vy, vz,
                            rho_in, bx, by, bz, psi, gam, kern);
    }
}

template <int CH, bool MIXED, bool MHD, int KT>
__global__ __launch_bounds__(64) void k_forces_cell(const RunDev* __restrict__ runs, int nruns, int ncell_tot,
                                                    int nblk_total, StateBuf S, WorkBuf W, int kmax)
{
    constexpr int NP = MHD ? 5 : 4;
    constexpr int RMAX = 1, NS = 27;
    __shared__ NbListR<RMAX> L;
    __shared__ float4 sp[NP][CH];
    __shared__ uint16_t lst[kCellListCap];
    const int cell = xcd_cell(blockIdx.x, nblk_total, ncell_tot);
    if (cell < 0) return;
    const int t0 = (int)W.cs[cell];
    const int nt = (int)W.cs[cell + 1] - t0;
    if (nt == 0) return;
    const int r = find_run_by_cell(runs, nruns, cell);
    const RunDev& R = runs[r];
    build_nblist<MIXED, RMAX>(R, cell - R.c0, W.cs, L);
    const int total = L.off[NS];
    const int lane = threadIdx.x;
    const LaneMap lm = lane_map(nt, kmax);
    const int K = lm.K, k = lm.k;
    const int Lr = (total + K - 1) / K;
    const bool keep = nt <= lm.tpg && Lr <= 32 * kMaskWords;  // the density pass stored masks
    const int kern = KT >= 0 ? KT : R.kernel;  // compile-time when every run uses one kernel
    const float H2 = R.H2;
    const float alpha = R.alpha, beta = R.beta;
    const bool mhd = MHD && R.mhd;
    const bool clean = mhd && R.cleaning;
    const float borve = mhd ? R.borve : 0.0f;
    const float ch = __uint_as_float(W.chbits[r]);
    const uint32_t* mrow = W.nbmask + (size_t)cell * (kMaskWords * 64) + lane;

    for (int tg = 0; tg < nt; tg += lm.tpg) {
        const int t = tg + lm.tl;
        const bool valid = lm.inwave && t < nt;
        const int i = t0 + (valid ? t : 0);
        const bool tgt = valid && is_target(S.species[i], S.mask[i]);
        const float4 o0 = W.P0[i], o1 = W.P1[i], o2 = W.P2[i], o3 = W.P3[i];
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MHD) o4 = W.P4[i];
        const int self_v = L.off[self_segment<RMAX>(R)] + t;
        const int v0 = lm.inwave ? k * Lr : total;
        const int v1 = v0 + Lr < total ? v0 + Lr : total;
        ForceAcc A = {};

        // neighbour lists from the stored masks: lane order = candidate order within a target
        const int mine = (keep && tgt) ? mask_count(mrow, 0, v1 - v0) : 0;
        const int incl = wave_incl_scan(mine);
        const bool listed = keep && __shfl(incl, 63, 64) <= kCellListCap;
        int lo = 0;
        if (listed) {
            int pos = incl - mine;
            for (int w = 0; w * 32 < v1 - v0 && mine > 0; ++w) {
                const int lim = (v1 - v0) - w * 32 < 32 ? (v1 - v0) - w * 32 : 32;
                uint32_t m = mask_bits(mrow, w * 32, lim);
                while (m) {
                    const int ii = __builtin_ctz(m);
                    m &= m - 1;
                    lst[pos++] = (uint16_t)(v0 + w * 32 + ii);
                }
            }
            // every lane takes part in the shuffle (a lane reading from an inactive lane does not
            // get that lane's value): clamp the source, then select
            const int pre = __shfl(incl, lm.gfirst > 0 ? lm.gfirst - 1 : 0, 64);
            lo = lm.gfirst > 0 ? pre : 0;  // first entry of this target's list
        }

        for (int base = 0; base < total; base += CH) {
            const int top = base + CH < total ? base + CH : total;
            for (int v = base + lane; v < top; v += 64) {
                const int e = v - base;
                const int sg = nb_segment(L, v);
                const int j = L.start[sg] + (v - L.off[sg]);
                float4 p = W.P0[j];
                p.x += L.sh[sg][0];
                p.y += L.sh[sg][1];
                p.z += L.sh[sg][2];
                sp[0][e] = p;
                sp[1][e] = W.P1[j];
                sp[2][e] = W.P2[j];
                sp[3][e] = W.P3[j];
                if constexpr (MHD) sp[4][e] = W.P4[j];
            }
            __syncthreads();
            if (listed) {
                // this target's entries below `top`: counted from each sub-lane's own bits
                const int a = v0 > base ? v0 : base;
                const int b = v1 < top ? v1 : top;
                const int in_chunk = (tgt && b > a) ? mask_count(mrow, a - v0, b - v0) : 0;
                const int ci = wave_incl_scan(in_chunk);
                const int cpre = __shfl(ci, lm.gfirst > 0 ? lm.gfirst - 1 : 0, 64);
                const int clast = __shfl(ci, lm.glast, 64);
                const int g = lm.inwave ? clast - (lm.gfirst > 0 ? cpre : 0) : 0;
                for (int q = k; q < g; q += K) {
                    const int e = (int)lst[lo + q] - base;
                    float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                    if constexpr (MHD) q4 = sp[4][e];
                    force_pair<MHD>(o0, o1, o2, o3, o4, sp[0][e], sp[1][e], sp[2][e], sp[3][e], q4, kern, alpha,
                                    beta, mhd, clean, A);
                }
                lo += g;
            } else {
                const int a = v0 > base ? v0 : base;
                const int b = v1 < top ? v1 : top;
                for (int cc = a; cc < b; cc += 32) {
                    const int lim = b - cc < 32 ? b - cc : 32;
                    uint32_t msk;
                    if (keep) {
                        msk = mask_bits(mrow, cc - v0, lim);
                    } else {
                        msk = cull32(sp[0], cc - base, 1, lim, o0.x, o0.y, o0.z, H2);
                        if (self_v >= cc && self_v < cc + lim) msk &= ~(1u << (self_v - cc));
                    }
                    if (!tgt) msk = 0;
                    while (msk) {
                        const int ii = __builtin_ctz(msk);
                        msk &= msk - 1;
                        const int e = cc - base + ii;
                        float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                        if constexpr (MHD) q4 = sp[4][e];
                        force_pair<MHD>(o0, o1, o2, o3, o4, sp[0][e], sp[1][e], sp[2][e], sp[3][e], q4, kern, alpha,
                                        beta, mhd, clean, A);
                    }
                }
            }
            __syncthreads();
        }
        force_finish<MHD>(R, W, lm, A, tgt, i, o1, o2, o3, o4, mhd, clean, borve, ch);
    }
}

// ------------------------------------------------------------------------------------------------
// kick + CFL + FP64 conservation sums (deterministic two-level reduction)
// ------------------------------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(256) void k_kick(const RunDev* __restrict__ runs, StateBuf S, WorkBuf W, double dt_host,
                                              int do_kick)
{
    __shared__ double red[6][4];
    const int b = blockIdx.x;
    const int r = W.blk_run[b];
    const RunDev& R = runs[r];
    const int64_t local = (int64_t)(b - R.kb_first) * 256 + threadIdx.x;
    double v_dt = INFINITY, v_m = 0, v_px = 0, v_py = 0, v_pz = 0, v_e = 0;
    if (local < R.np) {
        const int64_t i = R.p0 + local;
        if (is_target(S.species[i], S.mask[i])) {
            const double dt = dt_host > 0.0 ? dt_host : W.dt[r];
            const float dtf = (float)dt;
            float vx = S.vel[0][i], vy = S.vel[1][i], vz = S.vel[2][i];
            float u = S.u[i];
            const float ax = W.dvdt[0][i], ay = W.dvdt[1][i], az = W.dvdt[2][i];
            float bx = 0.f, by = 0.f, bz = 0.f;
            const bool kick = do_kick && !W.prime_flag[r] && isfinite(dt);  // no CFL bound: no kick
            if (kick) {
                vx = vx + ax * dtf;
                vy = vy + ay * dtf;
                vz = vz + az * dtf;
                u = u + W.dudt[i] * dtf;
                S.vel[0][i] = vx;
                S.vel[1][i] = vy;
                S.vel[2][i] = vz;
                S.u[i] = u;
            }
            if (R.mhd) {
                bx = S.B[0][i];
                by = S.B[1][i];
                bz = S.B[2][i];
                if (kick) {
                    bx = bx + W.dbdt[0][i] * dtf;
                    by = by + W.dbdt[1][i] * dtf;
                    bz = bz + W.dbdt[2][i] * dtf;
                    S.B[0][i] = bx;
                    S.B[1][i] = by;
                    S.B[2][i] = bz;
                    S.psi[i] = S.psi[i] + W.dpsidt[i] * dtf;
                }
            }
            const float h = S.h[i], vsg = W.vsig[i];
            if (MIXED) {
                double dti = (double)R.cfl * (double)h / (double)vsg;
                double a2 = (double)ax * ax + (double)ay * ay + (double)az * az;
                if (a2 > 0.0) dti = fmin(dti, (double)R.cfl_force * sqrt((double)h / sqrt(a2)));
                v_dt = dti;
            } else {
                float dti = R.cfl * h / vsg;
                float a2 = ax * ax + ay * ay + az * az;
                if (a2 > 0.0f) dti = fminf(dti, R.cfl_force * sqrtf(h / sqrtf(a2)));
                v_dt = (double)dti;
            }
            const double m = (double)S.mass[i];
            const double dvx = vx, dvy = vy, dvz = vz;
            const double b2 = (double)bx * bx + (double)by * by + (double)bz * bz;
            v_m = m;
            v_px = m * dvx;
            v_py = m * dvy;
            v_pz = m * dvz;
            v_e = m * (0.5 * (dvx * dvx + dvy * dvy + dvz * dvz) + (double)u + 0.5 * b2 / (double)S.rho[i]);
        }
    }
    // wave then block reduction, fixed order
    for (int off = 32; off >= 1; off >>= 1) {
        v_dt = fmin(v_dt, __shfl_xor(v_dt, off, 64));
        v_m += __shfl_xor(v_m, off, 64);
        v_px += __shfl_xor(v_px, off, 64);
        v_py += __shfl_xor(v_py, off, 64);
        v_pz += __shfl_xor(v_pz, off, 64);
        v_e += __shfl_xor(v_e, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = v_dt;
        red[1][wid] = v_m;
        red[2][wid] = v_px;
        red[3][wid] = v_py;
        red[4][wid] = v_pz;
        red[5][wid] = v_e;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        double s;
        if (q == 0) s = fmin(fmin(red[0][0], red[0][1]), fmin(red[0][2], red[0][3]));
        else s = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
        W.partial[(int64_t)b * 6 + q] = s;
    }
}

// One block per run: reduce the run's kick-block partials in a fixed tree order, publish the
// conserved sums, set the next dt and advance the time.
__global__ __launch_bounds__(256) void k_finalize(const RunDev* __restrict__ runs, WorkBuf W, double dt_host,
                                                  int advance)
{
    __shared__ double red[6][4];
    const int r = blockIdx.x;
    const RunDev& R = runs[r];
    double s[6] = {INFINITY, 0, 0, 0, 0, 0};
    const double* base = W.partial + (int64_t)R.kb_first * 6;
    int bb = threadIdx.x;
    for (; bb + 768 < R.kb_count; bb += 1024) {  // four partials per round, loads independent
        double v[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 6; ++q) v[u][q] = base[(int64_t)(bb + 256 * u) * 6 + q];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s[0] = fmin(s[0], v[u][0]);
            for (int q = 1; q < 6; ++q) s[q] += v[u][q];
        }
    }
    for (; bb < R.kb_count; bb += 256) {
        const double* p = base + (int64_t)bb * 6;
        s[0] = fmin(s[0], p[0]);
        for (int q = 1; q < 6; ++q) s[q] += p[q];
    }
    // butterfly within each wave (every lane ends with the wave's result), then the four waves in order
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        s[0] = fmin(s[0], __shfl_xor(s[0], off, 64));
#pragma unroll
        for (int q = 1; q < 6; ++q) s[q] += __shfl_xor(s[q], off, 64);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 6; ++q) red[q][wid] = s[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            red[0][0] = fmin(red[0][0], red[0][w]);
            for (int q = 1; q < 6; ++q) red[q][0] += red[q][w];
        }
        const double dt_used = dt_host > 0.0 ? dt_host : W.dt[r];
        const int priming = W.prime_flag[r];
        double t = W.time[r];
        if (advance && !priming && isfinite(dt_used)) t += dt_used;
        W.time[r] = t;
        const double dtn = R.dt_fixed > 0.0 ? R.dt_fixed : red[0][0];
        double* c = W.cons + (int64_t)r * 8;
        c[0] = red[1][0];
        c[1] = red[2][0];
        c[2] = red[3][0];
        c[3] = red[4][0];
        c[4] = red[5][0];
        c[5] = dtn;
        c[6] = t;
        c[7] = (double)__uint_as_float(W.chbits[r]);
        // a step sets the next dt; a reduction outside a step (prime, queries) sets it only for
        // runs still waiting for their first dt, so runs already stepping keep their own sequence.
        // A run's first step after new initial data is its priming step: no drift, no kick, no
        // time advance (the same arithmetic as sph_prime), only its first CFL dt
        if (advance || priming) W.dt[r] = dtn;
        if (advance) W.prime_flag[r] = 0;  // the step that primed the run is done
    }
}

// Pool compaction: destination particle d belongs to the last kept run whose destination start is
// <= d; every state column is copied from its source slot.
__global__ __launch_bounds__(256) void k_move_runs(StateBuf A, StateBuf B, const int64_t* __restrict__ tab, int n_keep,
                                                   int64_t total, int mixed, int mhd)
{
    const int64_t d = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (d >= total) return;
    int lo = 0, hi = n_keep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[3 * mid] <= d) lo = mid;
        else hi = mid - 1;
    }
    const int64_t s = tab[3 * lo + 1] + (d - tab[3 * lo]);
    B.id[d] = A.id[s];
    B.mask[d] = A.mask[s];
    B.species[d] = A.species[s];
    for (int a = 0; a < 3; ++a) {
        if (mixed) B.pos64[a][d] = A.pos64[a][s];
        else B.pos32[a][d] = A.pos32[a][s];
        B.vel[a][d] = A.vel[a][s];
        if (mhd) B.B[a][d] = A.B[a][s];
    }
    B.mass[d] = A.mass[s];
    B.h[d] = A.h[s];
    B.u[d] = A.u[s];
    B.rho[d] = A.rho[s];
    if (mhd) B.psi[d] = A.psi[s];
}

// Per-run scalars across a change of the run set: the old runs' dt, time and priming flag are saved,
// then new run r takes those of old run run_src[r], or starts fresh when run_src[r] < 0 (dt = its
// fixed dt or 0 until its priming step sets one, time 0, primed by its next step).
__global__ __launch_bounds__(256) void k_save_run_scalars(WorkBuf W, int n_old)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n_old) return;
    W.run_tmp[3 * q] = W.dt[q];
    W.run_tmp[3 * q + 1] = W.time[q];
    W.run_tmp[3 * q + 2] = (double)W.prime_flag[q];
}

__global__ __launch_bounds__(256) void k_load_run_scalars(const RunDev* __restrict__ runs, WorkBuf W, int n_new)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n_new) return;
    const int q = W.run_src[r];
    if (q >= 0) {
        W.dt[r] = W.run_tmp[3 * q];
        W.time[r] = W.run_tmp[3 * q + 1];
        W.prime_flag[r] = (int32_t)W.run_tmp[3 * q + 2];
    } else {
        W.dt[r] = runs[r].dt_fixed > 0.0 ? runs[r].dt_fixed : 0.0;
        W.time[r] = 0.0;
        W.prime_flag[r] = 1;
    }
}

// New initial data for one run: time 0, dt its fixed dt or 0, primed by its next step.
__global__ void k_restart_run(const RunDev* __restrict__ runs, WorkBuf W, int run)
{
    W.dt[run] = runs[run].dt_fixed > 0.0 ? runs[run].dt_fixed : 0.0;
    W.time[run] = 0.0;
    W.prime_flag[run] = 1;
}

hipError_t launch_remap_runs(const RunDev* runs, const WorkBuf& W, int n_old, int n_new, hipStream_t s)
{
    if (n_old > 0) k_save_run_scalars<<<(unsigned)((n_old + 255) / 256), 256, 0, s>>>(W, n_old);
    if (n_new > 0) k_load_run_scalars<<<(unsigned)((n_new + 255) / 256), 256, 0, s>>>(runs, W, n_new);
    return hipGetLastError();
}

hipError_t launch_restart_run(const RunDev* runs, const WorkBuf& W, int run, hipStream_t s)
{
    k_restart_run<<<1, 1, 0, s>>>(runs, W, run);
    return hipGetLastError();
}

hipError_t launch_move_runs(const StateBuf& A, const StateBuf& B, const int64_t* tab, int n_keep, int64_t total,
                            int mixed, int mhd, hipStream_t s)
{
    if (total <= 0 || n_keep <= 0) return hipSuccess;
    k_move_runs<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(A, B, tab, n_keep, total, mixed, mhd);
    return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline int nblk_xcd(int ncell) { return ((ncell + 7) / 8) * 8; }

// Blocks of a chunked pair-pass launch: one wave per supercell (persistent_waves 0), or a persistent
// grid of that many waves (a multiple of 8) striding over the supercells when there are more cells
// than waves. The whole-cell passes and the 125-cell stencil launch one wave per supercell.
static inline int pair_grid(const LaunchCtx& c, int nblk)
{
    return (c.persistent_waves > 0 && c.persistent_waves < nblk) ? c.persistent_waves : nblk;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_drift_bin(const LaunchCtx& c, double dt_host, int do_drift)
{
    if (c.mixed)
        k_drift_bin<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_drift);
    else
        k_drift_bin<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_drift);
    return hipGetLastError();
}

hipError_t launch_scan(const LaunchCtx& c)
{
    const int n = c.ncell_tot;
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    k_scan_local<<<nb, 256, 0, c.stream>>>(c.W.count, c.W.cs, c.W.bsum, n);
    k_scan_top<<<1, 1024, 0, c.stream>>>(c.W.bsum, nb, c.W.cs, n);
    k_scan_add<<<nb, 256, 0, c.stream>>>(c.W.cs, c.W.bsum, n);
    return hipGetLastError();
}

hipError_t launch_scatter_gather(const LaunchCtx& c, const StateBuf& Snew)
{
    k_scatter<<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W);
    if (c.mixed)
        k_gather<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, Snew, c.W, c.has_mhd);
    else
        k_gather<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, Snew, c.W, c.has_mhd);
    return hipGetLastError();
}

#define SPH_COMMA ,
// Instantiate a pair kernel for (precision, MHD fields, smoothing kernel): KT = -1 reads the kernel
// type per run (batches that mix kernel types).
#define SPH_PAIR_DISPATCH(KERNEL, GRIDN, SIZE, EXTRA, ...)                                          \
    do {                                                                                            \
        auto go = [&](auto mixed, auto mhd, auto kt) {                                              \
            KERNEL<SIZE, decltype(mixed)::value, decltype(mhd)::value, decltype(kt)::value EXTRA>    \
                <<<(GRIDN), 64, 0, c.stream>>>(__VA_ARGS__);                                         \
        };                                                                                          \
        auto kt_go = [&](auto mixed, auto mhd) {                                                    \
            switch (c.kern) {                                                                       \
            case 0: go(mixed, mhd, std::integral_constant<int, 0>{}); break;                        \
            case 1: go(mixed, mhd, std::integral_constant<int, 1>{}); break;                        \
            case 2: go(mixed, mhd, std::integral_constant<int, 2>{}); break;                        \
            default: go(mixed, mhd, std::integral_constant<int, -1>{}); break;                      \
            }                                                                                       \
        };                                                                                          \
        if (c.mixed) {                                                                              \
            if (c.has_mhd) kt_go(std::true_type{}, std::true_type{});                               \
            else kt_go(std::true_type{}, std::false_type{});                                        \
        } else {                                                                                    \
            if (c.has_mhd) kt_go(std::false_type{}, std::true_type{});                              \
            else kt_go(std::false_type{}, std::false_type{});                                       \
        }                                                                                           \
    } while (0)

template <int CH>
static void density_ch(const LaunchCtx& c, int nblk)
{
    const int grid = pair_grid(c, nblk);
    if (grid < nblk)
        SPH_PAIR_DISPATCH(k_density, grid, CH, SPH_COMMA 1 SPH_COMMA true, c.runs, c.nruns, c.ncell_tot, nblk, c.S,
                          c.W, c.kmax);
    else
        SPH_PAIR_DISPATCH(k_density, nblk, CH, , c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax);
}

// supercells finer than H somewhere in the context: the chunked lane loop over the 125-cell stencil
template <int CH>
static void density_fine_ch(const LaunchCtx& c, int nblk)
{
    SPH_PAIR_DISPATCH(k_density, nblk, CH, SPH_COMMA 2, c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax);
}

template <int CAPD>
static void density_cell_ch(const LaunchCtx& c, int nblk)
{
    SPH_PAIR_DISPATCH(k_density_cell, nblk, CAPD, , c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax);
}

template <int CH>
static void forces_cell_ch(const LaunchCtx& c, int nblk)
{
    SPH_PAIR_DISPATCH(k_forces_cell, nblk, CH, , c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax);
}

hipError_t launch_density(const LaunchCtx& c)
{
    const int nblk = nblk_xcd(c.ncell_tot);
    if (c.rmax == 2) {
        if (c.density_chunk >= 256) density_fine_ch<256>(c, nblk);
        else density_fine_ch<128>(c, nblk);
    } else if (c.balance == 2) {
        density_cell_ch<512>(c, nblk);
    } else if (c.density_chunk >= 256) {
        density_ch<256>(c, nblk);
    } else {
        density_ch<128>(c, nblk);
    }
    return hipGetLastError();
}

template <int CH, int BAL>
static void forces_ch(const LaunchCtx& c, int nblk)
{
    const int grid = pair_grid(c, nblk);
    if (grid < nblk)
        SPH_PAIR_DISPATCH(k_forces, grid, CH, SPH_COMMA BAL SPH_COMMA 1 SPH_COMMA true, c.runs, c.nruns, c.ncell_tot,
                          nblk, c.S, c.W, c.kmax, c.balance);
    else if (c.one_h)  // one smoothing length per run: the pair loop neither tests nor averages h
        SPH_PAIR_DISPATCH(k_forces, nblk, CH, SPH_COMMA BAL SPH_COMMA 1 SPH_COMMA false SPH_COMMA true, c.runs, c.nruns,
                          c.ncell_tot, nblk, c.S, c.W, c.kmax, c.balance);
    else
        SPH_PAIR_DISPATCH(k_forces, nblk, CH, SPH_COMMA BAL, c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W, c.kmax,
                          c.balance);
}

template <int CH>
static void forces_fine_ch(const LaunchCtx& c, int nblk)
{
    SPH_PAIR_DISPATCH(k_forces, nblk, CH, SPH_COMMA 0 SPH_COMMA 2, c.runs, c.nruns, c.ncell_tot, nblk, c.S, c.W,
                      c.kmax, 0);
}

hipError_t launch_forces(const LaunchCtx& c)
{
    const int nblk = nblk_xcd(c.ncell_tot);
    if (c.rmax == 2) {
        if (c.force_chunk >= 128) forces_fine_ch<128>(c, nblk);
        else forces_fine_ch<64>(c, nblk);
    } else if (c.balance == 2) {
        forces_cell_ch<128>(c, nblk);
    } else if (c.force_chunk >= 128) {
        if (c.balance == 1) forces_ch<128, 1>(c, nblk);
        else forces_ch<128, 0>(c, nblk);
    } else {
        if (c.balance == 1) forces_ch<64, 1>(c, nblk);
        else forces_ch<64, 0>(c, nblk);
    }
    return hipGetLastError();
}

hipError_t launch_kick(const LaunchCtx& c, double dt_host, int do_kick)
{
    if (c.mixed)
        k_kick<true><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_kick);
    else
        k_kick<false><<<c.nkb, 256, 0, c.stream>>>(c.runs, c.S, c.W, dt_host, do_kick);
    k_finalize<<<c.nruns, 256, 0, c.stream>>>(c.runs, c.W, dt_host, do_kick);
    return hipGetLastError();
}

}  // namespace sph
